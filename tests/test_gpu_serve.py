"""The request server (coupe_hip_server_*, coupe_hip_mesh_*): every result
equals the oracle's ids on the same input, alone or in the mix; resident
meshes repartitioned after drift; deep requests among shallow ones; poll,
cancel, double waits and concurrent submitters; admission under a budget.

Integer-valued weights throughout: ids must equal the oracle's bit for bit.
"""
import sys
import threading

import numpy as np
import pytest

import coupe_amd
from coupe_amd import _native, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def oracle_ids(oracle, pts, w, iters, tol):
    err, ids, _ = oracle.rcb(pts, w, iters, tol, parallel=pts.shape[0] > 200_000, nthreads=16)
    assert err == 0
    return ids


def as_u64(out):
    return out.cpu().numpy().astype(np.uint64) if hasattr(out, "cpu") else out


def test_server_results_equal_oracle(oracle_mod):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(21)
    reqs = []
    for q in range(60):
        n = int(10 ** rng.uniform(2, 5.3))
        dim = 3 if q % 4 else 2
        pts = rng.random((n, dim))
        w = rng.integers(1, 100, n).astype([np.float64, np.int32, np.int64][q % 3])
        reqs.append((pts, w, int(rng.integers(3, 11)), float(rng.choice([0.01, 0.05]))))
    say(f"server: {len(reqs)} requests")
    srv = coupe_amd.Server(slots=3, memory_budget=8 << 30, max_batch_requests=16)
    try:
        outs, tickets = [], []
        for q, (pts, w, it, tol) in enumerate(reqs):
            if q % 3 == 0:   # device tensors in and out
                out = torch.zeros(pts.shape[0], dtype=torch.int64, device="cuda")
                t = srv.submit(out, torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda(), it, tol)
            else:            # host buffers
                out = np.zeros(pts.shape[0], dtype=np.uint64)
                t = srv.submit(out, pts, w, it, tol)
            outs.append(out)
            tickets.append(t)
        stats = [srv.wait(t) for t in tickets]
        assert srv.stats()["completed"] == len(reqs)
    finally:
        srv.close()
    assert max(s["batch_requests"] for s in stats) > 1   # requests were batched
    for q, ((pts, w, it, tol), out) in enumerate(zip(reqs, outs)):
        assert np.array_equal(as_u64(out), oracle_ids(oracle_mod, pts, w, it, tol)), f"request {q}"


def test_served_stream_with_resident_meshes(oracle_mod):
    """The served workload's generator (sizes log-uniform in [10^4, 10^6]
    here, 3D points uniform in [0,1)^3, weights 1..100, iter_count 6..12,
    tolerance 0.01): new requests from host memory mixed with repartitions
    of resident meshes, each after every point moved by N(0, 0.01) and 5% of
    its weights were resampled (updates from host arrays and from GPU
    tensors).  Every result equals the oracle on the mesh's current points;
    the mesh's grouping by part is the stable grouping of its ids."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    srv = coupe_amd.Server(slots=4, memory_budget=16 << 30, max_batch_requests=32)
    meshes = []
    try:
        for k in range(4):
            n = int(round(10 ** rng.uniform(4, 6)))
            p, w = workloads.uniform(n, 3, 4000 + k, 100)
            meshes.append([srv.create_mesh(p, w) if k % 2 else
                           srv.create_mesh(torch.from_numpy(p).cuda(), torch.from_numpy(w).cuda()), p, w])
        checks = []
        for r in range(40):
            iters = int(6 + np.floor(7 * rng.random()))
            if rng.random() < 0.3:
                k = int(rng.integers(0, len(meshes)))
                mesh, p, w = meshes[k]
                p = p + 0.01 * rng.standard_normal(p.shape)
                w = w.copy()
                sel = rng.random(w.shape[0]) < 0.05
                w[sel] = np.floor(rng.random(int(sel.sum())) * 100) + 1
                if r % 2:
                    mesh.update(points=p, weights=w)
                else:
                    mesh.update(points=torch.from_numpy(p).cuda(), weights=torch.from_numpy(w).cuda())
                meshes[k][1:] = [p, w]
                out = (np.zeros(p.shape[0], dtype=np.uint64) if r % 3 else
                       torch.zeros(p.shape[0], dtype=torch.int64, device="cuda"))
                say(f"served stream {r}: repartition of mesh {k}, n={p.shape[0]} L={iters}")
                t = srv.submit_mesh(mesh, out, iters, 0.01)
                checks.append((t, out, p, w, iters, mesh))
            else:
                n = int(round(10 ** rng.uniform(4, 6)))
                p, w = workloads.uniform(n, 3, 5000 + r, 100)
                out = np.zeros(n, dtype=np.uint64)
                say(f"served stream {r}: new request, n={n} L={iters}")
                t = srv.submit(out, p, w, iters, 0.01)
                checks.append((t, out, p, w, iters, None))
        for t, out, p, w, iters, mesh in checks:
            srv.wait(t)
            exp = oracle_ids(oracle_mod, p, w, iters, 0.01)
            assert np.array_equal(as_u64(out), exp), f"ticket {t}: ids differ from the oracle"
        for mesh, p, w in meshes:
            # the last partition of the mesh, grouped by part
            t = srv.submit_mesh(mesh, None, 9, 0.01)
            srv.wait(t)
            perm, off = mesh.parts()
            ids = oracle_ids(oracle_mod, p, w, 9, 0.01)
            assert np.array_equal(perm, np.argsort(ids, kind="stable").astype(np.uint32))
            assert np.array_equal(off, np.searchsorted(np.sort(ids), np.arange(int(ids.max()) + 2)).astype(np.uint64))
            mesh.destroy()
        st = srv.stats()
        assert st["meshes"] == 0 and st["peak_device_bytes"] <= st["memory_budget"]
    finally:
        srv.close()


def test_deep_request_among_shallow_ones(oracle_mod):
    """A request with iter_count > 24 runs alone; its neighbours still batch."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(31)
    srv = coupe_amd.Server(slots=2, memory_budget=8 << 30, max_batch_requests=16)
    try:
        reqs = []
        for q in range(12):
            n = int(rng.integers(2_000, 60_000))
            pts = rng.random((n, 3))
            w = rng.integers(1, 50, n).astype(np.int64)
            iters = 27 if q == 5 else int(rng.integers(3, 9))
            out = np.zeros(n, dtype=np.uint64)
            reqs.append((srv.submit(out, pts, w, iters, 0.02), out, pts, w, iters))
        for t, out, pts, w, iters in reqs:
            srv.wait(t)
            assert np.array_equal(out, oracle_ids(oracle_mod, pts, w, iters, 0.02)), f"iter_count {iters}"
    finally:
        srv.close()


def test_poll_cancel_and_double_wait():
    rng = np.random.default_rng(41)
    srv = coupe_amd.Server(slots=1, memory_budget=4 << 30, max_batch_requests=1)
    try:
        big = rng.random((2_000_000, 3))
        wb = np.ones(2_000_000)
        out_big = np.zeros(2_000_000, dtype=np.uint64)
        t_big = srv.submit(out_big, big, wb, 12, 0.01)
        small = []
        for q in range(20):
            p = rng.random((1000, 2))
            out = np.zeros(1000, dtype=np.uint64)
            small.append((srv.submit(out, p, np.ones(1000), 4, 0.05), out, p))
        cancelled = [t for t, _, _ in small[10:] if srv.cancel(t)]
        assert cancelled, "queued requests behind a running one can be withdrawn"
        for t in cancelled:
            with pytest.raises(coupe_amd.CoupeError) as e:
                srv.poll(t)
            assert e.value.code == _native.NOT_FOUND
        srv.wait(t_big)
        for t, out, p in small:
            if t in cancelled:
                continue
            srv.wait(t)
            with pytest.raises(coupe_amd.CoupeError) as e:   # waited for: unknown
                srv.cancel(t)
            assert e.value.code == _native.NOT_FOUND
        with pytest.raises(coupe_amd.CoupeError) as e:
            srv.wait(t_big)
        assert e.value.code == _native.NOT_FOUND
        st = srv.stats()
        assert st["cancelled"] == len(cancelled) and st["queued"] == 0
    finally:
        srv.close()


def test_concurrent_submitters_and_waiters(oracle_mod):
    """Eight threads submit and wait at once (the ticket table is shared)."""
    rng = np.random.default_rng(51)
    srv = coupe_amd.Server(slots=3, memory_budget=8 << 30, max_batch_requests=8)
    cases = []
    for q in range(160):
        n = int(rng.integers(1, 4000))
        cases.append((rng.random((n, 2 + q % 2)), rng.integers(1, 9, n).astype(np.int64), int(rng.integers(1, 8))))
    results = {}
    errors = []

    def run(lo):
        try:
            mine = []
            for q in range(lo, len(cases), 8):
                pts, w, it = cases[q]
                out = np.zeros(pts.shape[0], dtype=np.uint64)
                mine.append((q, srv.submit(out, pts, w, it, 0.02), out))
            for q, t, out in mine:
                srv.wait(t)
                results[q] = out
        except Exception as e:   # reported by the main thread
            errors.append(e)

    try:
        ths = [threading.Thread(target=run, args=(k,)) for k in range(8)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
    finally:
        srv.close()
    assert not errors, errors
    for q, (pts, w, it) in enumerate(cases):
        assert np.array_equal(results[q], oracle_ids(oracle_mod, pts, w, it, 0.02)), f"case {q}"


def test_admission_under_a_small_budget(oracle_mod):
    """A budget that holds about one large request at a time: requests wait
    for memory, workers give theirs back, and every result is exact."""
    rng = np.random.default_rng(61)
    budget = 400 << 20
    srv = coupe_amd.Server(slots=3, memory_budget=budget, max_batch_requests=4)
    try:
        reqs = []
        for q in range(9):
            n = int(rng.integers(1_000_000, 2_500_000))
            pts = rng.random((n, 3))
            w = rng.integers(1, 100, n).astype(np.float64)
            out = np.zeros(n, dtype=np.uint64)
            reqs.append((srv.submit(out, pts, w, 10, 0.01), out, pts, w))
        for t, out, pts, w in reqs:
            srv.wait(t)
            assert np.array_equal(out, oracle_ids(oracle_mod, pts, w, 10, 0.01))
        st = srv.stats()
        assert st["peak_device_bytes"] <= budget, st
        with pytest.raises(coupe_amd.CoupeError) as e:   # cannot fit alone
            big = np.zeros((20_000_000, 3))
            srv.submit(np.zeros(20_000_000, dtype=np.uint64), big, np.ones(20_000_000), 4, 0.05)
        assert e.value.code == _native.ALLOC
    finally:
        srv.close()


def test_server_rejects_bad_requests():
    srv = coupe_amd.Server(slots=1, memory_budget=1 << 20)
    try:
        pts = np.zeros((100_000, 3))
        with pytest.raises(coupe_amd.CoupeError) as e:
            srv.submit(np.zeros(100_000, dtype=np.uint64), pts, np.ones(100_000), 4, 0.05)
        assert e.value.code == _native.ALLOC
        with pytest.raises(coupe_amd.CoupeError) as e:
            srv.submit(np.zeros(10, dtype=np.uint64), np.zeros((10, 4)), np.ones(10), 4, 0.05)
        assert e.value.code == _native.BAD_DIMENSION
    finally:
        srv.close()
