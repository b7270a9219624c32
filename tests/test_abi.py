"""The C ABI library on the host: it loads, exports every symbol the headers
declare, and returns coupe's status codes on the error paths that stop
before any device work (coupe-ffi/src/lib.rs:272-304)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import coupe_amd
from coupe_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# coupe-ffi/src/lib.rs:69-110
STRERROR = {
    0: "success",
    1: "allocation failed",
    2: "coupe encountered a bug and crashed",
    3: "this algorithm does not support the given mesh dimension",
    4: "this algorithm does not support the given type",
    5: "this algorithm does not support k-way partitioning",
    6: "input iters (e.g. weights and points) don't have the same length",
    7: "no partition has been found for the given constraints",
    8: "this algorithm does not support negative values",
}


def _declared_functions():
    names = []
    for h in ("coupe.h", "coupe_hip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(coupe_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_native.LIB_PATH)
    declared = _declared_functions()
    assert "coupe_rcb" in declared and "coupe_hip_rcb_device" in declared
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(declared) == sorted(_native.EXPORTS)


def test_enum_values_match_coupe():
    text = open(os.path.join(ROOT, "include", "coupe.h")).read()
    body = re.search(r"enum coupe_err \{(.*?)\};", text, re.S).group(1)
    names = re.findall(r"(COUPE_ERR_\w+)", body)
    assert names == ["COUPE_ERR_OK", "COUPE_ERR_ALLOC", "COUPE_ERR_CRASH", "COUPE_ERR_BAD_DIMENSION",
                     "COUPE_ERR_BAD_TYPE", "COUPE_ERR_BIPART_ONLY", "COUPE_ERR_LEN_MISMATCH",
                     "COUPE_ERR_NOT_FOUND", "COUPE_ERR_NEG_VALUES"]
    body = re.search(r"enum coupe_type \{(.*?)\};", text, re.S).group(1)
    assert re.findall(r"(COUPE_\w+)", body) == ["COUPE_INT", "COUPE_INT64", "COUPE_DOUBLE"]


def test_strerror_strings():
    for code, msg in STRERROR.items():
        assert _native.strerror(code) == msg


def test_data_sets_allocate_and_free():
    L = _native.lib()
    buf = np.zeros(8)
    one = C.c_int(1)
    d1 = L.coupe_data_array(8, _native.COUPE_DOUBLE, buf.ctypes.data)
    d2 = L.coupe_data_constant(8, _native.COUPE_INT, C.addressof(one))
    cb = _native.ITH_FN(lambda ctx, i: buf.ctypes.data + 8 * i)
    d3 = L.coupe_data_fn(None, 8, _native.COUPE_DOUBLE, cb)
    assert d1 and d2 and d3
    for d in (d1, d2, d3):
        L.coupe_data_free(d)
    L.coupe_data_free(None)


def test_rcb_error_paths_without_device():
    L = _native.lib()
    pts = np.zeros((4, 2))
    w = np.ones(3)
    part = np.full(4, 77, dtype=np.uint64)
    dp = L.coupe_data_array(4, _native.COUPE_DOUBLE, pts.ctypes.data)
    dw = L.coupe_data_array(3, _native.COUPE_DOUBLE, w.ctypes.data)
    assert L.coupe_rcb(part.ctypes.data, 2, dp, dw, 2, 0.05) == _native.LEN_MISMATCH
    L.coupe_data_free(dw)
    w4 = np.ones(4)
    dw = L.coupe_data_array(4, _native.COUPE_DOUBLE, w4.ctypes.data)
    for dim in (0, 1, 4, 7):
        assert L.coupe_rcb(part.ctypes.data, dim, dp, dw, 2, 0.05) == _native.BAD_DIMENSION
    # dimension is checked after the lengths, as in coupe
    assert L.coupe_rcb(part.ctypes.data, 5, dp, L.coupe_data_array(1, 2, w4.ctypes.data), 2, 0.05) \
        == _native.LEN_MISMATCH
    L.coupe_data_free(dp)
    L.coupe_data_free(dw)
    # empty input: OK, nothing written
    d0 = L.coupe_data_array(0, _native.COUPE_DOUBLE, pts.ctypes.data)
    w0 = L.coupe_data_array(0, _native.COUPE_DOUBLE, w4.ctypes.data)
    assert L.coupe_rcb(part.ctypes.data, 3, d0, w0, 5, 0.05) == _native.OK
    assert L.coupe_rcb(part.ctypes.data, 9, d0, w0, 5, 0.05) == _native.BAD_DIMENSION
    assert part.tolist() == [77] * 4
    L.coupe_data_free(d0)
    L.coupe_data_free(w0)


def test_python_mirror_errors():
    rcb = coupe_amd.Rcb(iter_count=2, tolerance=0.05)
    part = np.zeros(4, dtype=np.uint64)
    with pytest.raises(coupe_amd.InputLenMismatch) as e:
        rcb.partition(part, (np.zeros((4, 2)), np.ones(3)))
    assert (e.value.expected, e.value.actual) == (4, 3)
    with pytest.raises(coupe_amd.InputLenMismatch) as e:
        rcb.partition(part, (np.zeros((5, 2)), np.ones(4)))
    assert (e.value.expected, e.value.actual) == (4, 5)
    with pytest.raises(coupe_amd.CoupeError) as e:
        rcb.partition(part, (np.zeros((4, 4)), np.ones(4)))
    assert e.value.code == _native.BAD_DIMENSION
    assert rcb.partition(np.zeros(0, dtype=np.uint64), (np.zeros((0, 3)), np.ones(0))) is None


def test_workload_generator_is_splitmix64():
    wl = _native.workloads_lib()
    # SplitMix64 seeded with 0: published first outputs.
    assert [wl.wl_splitmix64(0, k) for k in range(3)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert wl.wl_uniform01(0, 0) == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53


def test_workloads_deterministic_and_in_range():
    from coupe_amd import workloads

    a, wa = workloads.gauss_mix(4096, seed=1)
    b, wb = workloads.gauss_mix(4096, seed=1)
    assert np.array_equal(a, b) and np.array_equal(wa, wb)
    assert wa.min() >= 1 and wa.max() <= 100 and np.all(wa == np.trunc(wa))
    g, wg = workloads.jgrid(16, seed=2)
    cell = np.floor(g * 16)
    idx = np.arange(16 ** 3)
    assert np.array_equal(cell[:, 0], idx % 16) and np.array_equal(cell[:, 2], idx // 256)
    assert np.all(np.abs(g * 16 - cell - 0.5) <= 0.25) and np.all(wg >= 1) and np.all(wg == np.trunc(wg))
    u, wu = workloads.uniform(1000, 2, 0)
    assert np.all((u >= 0) & (u < 1)) and np.all(wu == 1.0)
