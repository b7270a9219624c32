"""coupe_hip_rcb_device_noalloc: the entry point that allocates nothing
(every work buffer is a carve of one static device array per GPU) gives the
ids of coupe_hip_rcb_device, call after call on the shared array, up to its
stated limits, and rejects larger inputs."""
import ctypes as C
import sys

import numpy as np
import pytest

import coupe_amd
from coupe_amd import _native

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _inputs(rng, n, dim, kind):
    pts = rng.random((n, dim)) if kind != "gauss" else rng.normal(0, 1, (n, dim)) ** 3
    if kind == "real":
        w = rng.random(n) * 7.0 + 0.25
    elif kind == "wide":
        w = rng.integers(1, 1 << 40, n).astype(np.int64)
    elif kind == "int32":
        w = rng.integers(1, 1000, n).astype(np.int32)
    else:
        w = rng.integers(1, 100, n).astype(np.float64)
    return pts, w


def _noalloc(torch, pts_d, w_d, iters, tol, out=None):
    n, dim = pts_d.shape
    if out is None:
        out = torch.full((n,), -1, dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    wtype = coupe_amd._weight_type(w_d.dtype)
    err = _native.lib().coupe_hip_rcb_device_noalloc(C.c_void_p(st), out.data_ptr(), n, dim, pts_d.data_ptr(),
                                                      wtype, w_d.data_ptr(), iters, tol)
    return err, out


def _with_context(torch, ctx, pts_d, w_d, iters, tol):
    out = torch.full((pts_d.shape[0],), -1, dtype=torch.int64, device="cuda")
    coupe_amd.Rcb(iter_count=iters, tolerance=tol).partition(out, (pts_d, w_d), context=ctx)
    return out


# pairs of calls in a row, the second smaller than the first
SEQUENCES = [
    [(_native.NOALLOC_MAX_POINTS, 3, 12, "int"), (70_001, 3, 9, "int")],
    [(300_000, 2, 12, "real"), (4_097, 2, 5, "real")],
    [(_native.NOALLOC_MAX_POINTS, 3, 12, "int"), (_native.NOALLOC_MAX_POINTS - 5, 3, 12, "real")],
    [(_native.NOALLOC_MAX_POINTS, 3, 12, "wide"), (123_457, 3, 11, "gauss")],
    [(_native.NOALLOC_MAX_POINTS, 2, 12, "int32"), (1, 2, 3, "int32")],
]


@pytest.mark.parametrize("seq", range(len(SEQUENCES)))
def test_noalloc_equals_device_entry(seq):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(800 + seq)
    ctx = coupe_amd.DeviceContext(0)
    try:
        for k, (n, dim, iters, kind) in enumerate(SEQUENCES[seq]):
            say(f"no-alloc sequence {seq} call {k}: n={n} d={dim} L={iters} weights {kind}")
            pts, w = _inputs(rng, n, dim, kind)
            pts_d, w_d = torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()
            tol = 0.01
            err, got = _noalloc(torch, pts_d, w_d, iters, tol)
            assert err == _native.OK, _native.strerror(err)
            exp = _with_context(torch, ctx, pts_d, w_d, iters, tol)
            torch.cuda.synchronize()
            if kind != "real":
                assert torch.equal(got, exp), f"sequence {seq} call {k}: {int((got != exp).sum())} ids differ"
            elif not torch.equal(got, exp):
                # real weights are summed with atomics: per-part weights agree to 1e-12
                a = np.bincount(got.cpu().numpy(), weights=w)
                b = np.bincount(exp.cpu().numpy(), weights=w)
                assert a.shape == b.shape
                np.testing.assert_allclose(a, b, rtol=1e-12, atol=0)
    finally:
        ctx.close()


def test_noalloc_limits():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(99)
    n = _native.NOALLOC_MAX_POINTS + 1
    pts, w = _inputs(rng, n, 3, "int")
    pts_d, w_d = torch.from_numpy(pts).cuda(), torch.from_numpy(w).cuda()
    say("no-alloc limits: one point too many, one level too deep")
    err, out = _noalloc(torch, pts_d, w_d, 8, 0.01)
    assert err == _native.ALLOC
    err, out = _noalloc(torch, pts_d[:1000], w_d[:1000], _native.NOALLOC_MAX_ITER_COUNT + 1, 0.01)
    assert err == _native.ALLOC
    torch.cuda.synchronize()
    assert bool((out == -1).all()), "a rejected call wrote ids"
    # at the limits, and the trivial sizes
    err, out = _noalloc(torch, pts_d[:_native.NOALLOC_MAX_POINTS], w_d[:_native.NOALLOC_MAX_POINTS],
                        _native.NOALLOC_MAX_ITER_COUNT, 0.01)
    assert err == _native.OK
    err, _ = _noalloc(torch, pts_d[:0], w_d[:0], 5, 0.01)
    assert err == _native.OK
    err, out = _noalloc(torch, pts_d[:1000], w_d[:1000], 0, 0.01)
    assert err == _native.OK
    torch.cuda.synchronize()
    assert bool((out == 0).all())
