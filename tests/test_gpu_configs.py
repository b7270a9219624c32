"""Parity with the oracle on the benchmark workloads at their full sizes, and
on the paths only they reach: 4096 batched instances of 10^3..10^6 points,
the 2^27-point jittered grid, and trees deeper than the dense split heap
(iter_count 25..40, node-record descent).

Integer-valued weights throughout: ids must equal the oracle's bit for bit.
"""
import sys

import numpy as np
import pytest

import coupe_amd
from coupe_amd import workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def oracle_ids(oracle, pts, w, iters, tol):
    err, ids, _ = oracle.rcb(pts, w, iters, tol, parallel=True, nthreads=16)
    assert err == 0
    return ids


def batch_workload(count=4096, seed=3):
    """The batched workload: sizes log-uniform in [10^3, 10^6], D in {2, 3},
    a random box per instance, weights 1..16, iter_count 4..8.  Returns the
    instances in generator order with their points flat (each instance's
    n * D coordinates after the previous instance's)."""
    n, dim, lo, ext, iters = workloads.batch_header(count, seed=seed, lg_lo=3.0, lg_hi=6.0, l_lo=4, l_hi=8)
    pts, ws = [], []
    for q in range(count):
        d = int(dim[q])
        p, w = workloads.box(int(n[q]), d, lo[q, :d], ext[q, :d], workloads.instance_seed(seed, q), 16)
        pts.append(p.ravel())
        ws.append(w)
    return n, dim, iters, np.concatenate(pts), np.concatenate(ws)


@pytest.mark.slow
def test_batched_workload_full(oracle_mod):
    """The 4096 instances as one call of both dimensions, every instance
    against the oracle; then one call per dimension gives the same ids."""
    torch = pytest.importorskip("torch")
    tol = 0.02
    n, dim, iters, flat, w = batch_workload()
    specs = [(int(n[q]), int(dim[q]), int(iters[q]), tol) for q in range(len(n))]
    say(f"batched workload: {len(specs)} instances of D 2 and 3, {w.shape[0]} points, one call")
    out = torch.zeros(w.shape[0], dtype=torch.int64, device="cuda")
    d_flat, d_w = torch.from_numpy(flat).cuda(), torch.from_numpy(w).cuda()
    coupe_amd.partition_batch_mixed(out, d_flat, d_w, specs)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.uint64)
    off = poff = 0
    starts = []
    for q, (k, d, it, _) in enumerate(specs):
        starts.append((off, poff))
        p = flat[poff:poff + k * d].reshape(k, d)
        exp = oracle_ids(oracle_mod, p, w[off:off + k], it, tol)
        if not np.array_equal(got[off:off + k], exp):
            bad = int(np.count_nonzero(got[off:off + k] != exp))
            raise AssertionError(f"instance {q} (D={d}, n={k}, L={it}): {bad} ids differ")
        off += k
        poff += k * d
    say("batched workload: every instance equals the oracle")
    for d in (2, 3):
        idx = [q for q in range(len(specs)) if specs[q][1] == d]
        p = np.concatenate([flat[starts[q][1]:starts[q][1] + specs[q][0] * d] for q in idx]).reshape(-1, d)
        ww = np.concatenate([w[starts[q][0]:starts[q][0] + specs[q][0]] for q in idx])
        o = torch.zeros(ww.shape[0], dtype=torch.int64, device="cuda")
        coupe_amd.partition_batch(o, torch.from_numpy(p).cuda(), torch.from_numpy(ww).cuda(),
                                  [(specs[q][0], specs[q][2], tol) for q in idx])
        torch.cuda.synchronize()
        exp = np.concatenate([got[starts[q][0]:starts[q][0] + specs[q][0]] for q in idx])
        assert np.array_equal(o.cpu().numpy().astype(np.uint64), exp), f"one call per dimension, D={d}"


@pytest.mark.slow
def test_jgrid512_full(oracle_mod):
    pts, w, iters, tol = workloads.standard("jgrid512")
    say(f"jittered grid 512^3: n={pts.shape[0]} L={iters}")
    got = np.zeros(pts.shape[0], dtype=np.uint64)
    coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(got, (pts, w))
    exp = oracle_ids(oracle_mod, pts, w, iters, tol)
    assert np.array_equal(got, exp), f"{int(np.count_nonzero(got != exp))} ids differ"


DEEP = [  # (n, dim, iter_count, cloud, weights)
    (3_000, 2, 25, "uniform", "int32"),
    (50_000, 3, 26, "gauss", "int64"),
    (200_000, 2, 32, "grid", "float64"),
    (120_000, 3, 40, "uniform", "int64"),
    (7_777, 3, 33, "dup", "int32"),
]


@pytest.mark.parametrize("scan_direct", ["2048", "1"])
def test_deep_trees(oracle_mod, monkeypatch, scan_direct):
    """iter_count beyond the dense heap (24): split-tree node records and
    leaf ids.  COUPE_HIP_SCAN_DIRECT=1 makes every multi-block scan scan its
    block totals first (the path of very large part counts)."""
    torch = pytest.importorskip("torch")
    monkeypatch.setenv("COUPE_HIP_SCAN_DIRECT", scan_direct)
    ctx = coupe_amd.DeviceContext(0)
    try:
        for q, (n, dim, iters, cloud, wt) in enumerate(DEEP):
            rng = np.random.default_rng(4000 + q)
            if cloud == "uniform":
                pts = rng.random((n, dim))
            elif cloud == "gauss":
                pts = rng.normal(0, 1, (n, dim)) ** 3
            elif cloud == "grid":
                pts = np.round(rng.random((n, dim)) * 300) / 300
            else:
                pts = np.round(rng.random((n, dim)) * 8)   # many duplicates: leaves that cannot split
            w = rng.integers(1, 500, n).astype(wt)
            tol = float([0.0, 0.01, 0.05][q % 3])
            say(f"deep tree (scan_direct {scan_direct}) case {q}: n={n} d={dim} L={iters} {cloud} {wt}")
            out = torch.zeros(n, dtype=torch.int64, device="cuda")
            coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(
                out, (torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()), context=ctx)
            torch.cuda.synchronize()
            got = out.cpu().numpy().astype(np.uint64)
            err, exp, _ = oracle_mod.rcb(pts, w, iters, tol)
            assert err == 0
            assert np.array_equal(got, exp), f"case {q}: {int(np.count_nonzero(got != exp))} ids differ"
    finally:
        ctx.close()
