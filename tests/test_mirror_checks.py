"""The Python mirror's argument checks for GPU inputs, on the CPU.

A host pointer or a pointer of another device handed to a kernel would fault
the GPU, so the mirror raises TypeError before it takes any pointer, creates a
context or calls the library.  The tensors here are stand-ins that carry a
device, a dtype and a shape and fail the test if a pointer is taken.
"""
import numpy as np
import pytest
import torch

import coupe_amd


class FakeGpuTensor:
    is_cuda = True

    def __init__(self, shape, dtype, index=0, contiguous=True):
        self.shape = tuple(shape)
        self.dtype = dtype
        self.device = torch.device("cuda", index)
        self._contiguous = contiguous

    def __len__(self):
        return self.shape[0]

    def dim(self):
        return len(self.shape)

    @property
    def ndim(self):
        return len(self.shape)

    def numel(self):
        return int(np.prod(self.shape))

    def is_contiguous(self):
        return self._contiguous

    def contiguous(self):
        return self

    def data_ptr(self):
        raise AssertionError("a device pointer was taken before the argument checks")


class FakeContext:
    def __init__(self, device):
        self.device = device

    @property
    def handle(self):
        raise AssertionError("the context was used before the argument checks")


N = 16


def rcb(part_ids, points, weights, **kw):
    coupe_amd.Rcb(iter_count=2, tolerance=0.05).partition(part_ids, (points, weights), **kw)


def test_rcb_host_weights_with_device_points():
    with pytest.raises(TypeError, match="weights must be a GPU tensor"):
        rcb(FakeGpuTensor([N], torch.int64), FakeGpuTensor([N, 3], torch.float64), torch.ones(N, dtype=torch.int64))


def test_rcb_host_part_ids_with_device_points():
    with pytest.raises(TypeError):
        rcb(np.zeros(N, dtype=np.uint64), FakeGpuTensor([N, 3], torch.float64), FakeGpuTensor([N], torch.int64))


def test_rcb_tensors_on_two_devices():
    with pytest.raises(TypeError, match="several devices"):
        rcb(FakeGpuTensor([N], torch.int64), FakeGpuTensor([N, 3], torch.float64, index=1),
            FakeGpuTensor([N], torch.int64))


def test_rcb_part_ids_width_and_layout():
    for ids in (FakeGpuTensor([N], torch.int32), FakeGpuTensor([N], torch.int64, contiguous=False)):
        with pytest.raises(TypeError, match="64-bit"):
            rcb(ids, FakeGpuTensor([N, 3], torch.float64), FakeGpuTensor([N], torch.int64))


def test_rcb_context_on_another_device():
    with pytest.raises(TypeError, match="context"):
        rcb(FakeGpuTensor([N], torch.int64), FakeGpuTensor([N, 3], torch.float64), FakeGpuTensor([N], torch.int64),
            context=FakeContext(1))


def test_rcb_lengths_checked_first():
    with pytest.raises(coupe_amd.InputLenMismatch):
        rcb(FakeGpuTensor([N], torch.int64), FakeGpuTensor([N, 3], torch.float64), FakeGpuTensor([N - 1], torch.int64))


GRID = coupe_amd.Grid([4, 4])


def test_grid_partition_width_and_layout():
    for ids in (FakeGpuTensor([N], torch.int32), FakeGpuTensor([N], torch.int64, contiguous=False)):
        with pytest.raises(TypeError, match="64-bit"):
            GRID.rcb(ids, FakeGpuTensor([N], torch.float64), 2)
        with pytest.raises(TypeError, match="64-bit"):
            GRID.rcb(ids, np.ones(N), 2)


def test_grid_tensors_on_two_devices():
    with pytest.raises(TypeError, match="several devices"):
        GRID.rcb(FakeGpuTensor([N], torch.int64, index=1), FakeGpuTensor([N], torch.float64), 2)


def test_grid_context_on_another_device():
    with pytest.raises(TypeError, match="context"):
        GRID.rcb(FakeGpuTensor([N], torch.int64, index=1), np.ones(N), 2, context=FakeContext(0))
    with pytest.raises(TypeError, match="context"):
        GRID.rcb(np.zeros(N, dtype=np.uint64), FakeGpuTensor([N], torch.float64), 2, context=FakeContext(1))


def test_grid_cell_count():
    with pytest.raises(ValueError):
        GRID.rcb(FakeGpuTensor([N - 1], torch.int64), FakeGpuTensor([N], torch.float64), 2)


def test_batch_ids_width():
    with pytest.raises(TypeError, match="64-bit"):
        coupe_amd.partition_batch(FakeGpuTensor([N], torch.int32), FakeGpuTensor([N, 3], torch.float64),
                                  FakeGpuTensor([N], torch.int64), [(N, 2, 0.05)])
    with pytest.raises(TypeError, match="64-bit"):
        coupe_amd.partition_batch_mixed(FakeGpuTensor([N], torch.int32), FakeGpuTensor([3 * N], torch.float64),
                                        FakeGpuTensor([N], torch.int64), [(N, 3, 2, 0.05)])


def test_autotune_checks_like_the_batch():
    pts = FakeGpuTensor([3 * N], torch.float64)
    with pytest.raises(coupe_amd.InputLenMismatch):
        coupe_amd.autotune(FakeGpuTensor([N - 1], torch.int64), pts, FakeGpuTensor([N], torch.int64),
                           [(N, 3, 2, 0.05)])
    with pytest.raises(TypeError, match="coordinates"):
        coupe_amd.autotune(FakeGpuTensor([N], torch.int64), FakeGpuTensor([2 * N], torch.float64),
                           FakeGpuTensor([N], torch.int64), [(N, 3, 2, 0.05)])
    with pytest.raises(TypeError, match="context"):
        coupe_amd.autotune(FakeGpuTensor([N], torch.int64), pts, FakeGpuTensor([N], torch.int64),
                           [(N, 3, 2, 0.05)], context=FakeContext(1))


def server_stub(device=0):
    srv = object.__new__(coupe_amd.Server)   # the checks run before the server is used
    srv._s, srv._keep, srv.device = None, {}, device
    return srv


def test_server_submit_checks():
    srv = server_stub()
    pts, w = np.zeros((N, 3)), np.ones(N)
    for ids in (np.zeros(N, dtype=np.int32), np.zeros(2 * N, dtype=np.uint64)[::2],
                FakeGpuTensor([N], torch.int32), FakeGpuTensor([N], torch.int64, index=1)):
        with pytest.raises(TypeError):
            srv.submit(ids, pts, w, 2, 0.05)
    with pytest.raises(TypeError, match="float64"):
        srv.submit(np.zeros(N, dtype=np.uint64), FakeGpuTensor([N, 3], torch.float32),
                   FakeGpuTensor([N], torch.float64), 2, 0.05)
    with pytest.raises(TypeError, match="device"):
        server_stub(device=1).submit(np.zeros(N, dtype=np.uint64), FakeGpuTensor([N, 3], torch.float64),
                                     FakeGpuTensor([N], torch.float64), 2, 0.05)


def test_mesh_checks():
    srv = server_stub()
    with pytest.raises(TypeError, match="float64"):
        coupe_amd.Mesh(srv, FakeGpuTensor([N, 3], torch.float32), FakeGpuTensor([N], torch.float64))
    mesh = object.__new__(coupe_amd.Mesh)
    mesh.server, mesh.n, mesh.dimension, mesh.handle = srv, N, 3, 1
    mesh.weight_type = coupe_amd._native.COUPE_DOUBLE
    with pytest.raises(TypeError, match="float64"):
        mesh.update(points=FakeGpuTensor([4, 3], torch.float32))
    with pytest.raises(TypeError, match="device"):
        mesh.update(points=FakeGpuTensor([4, 3], torch.float64, index=1))
    with pytest.raises(TypeError, match="64-bit"):
        srv.submit_mesh(mesh, FakeGpuTensor([N], torch.int32), 2, 0.05)
