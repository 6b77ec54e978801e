"""Host-side tests of the drop-in boundary (no GPU compute).

* the C-ABI library loads and exports every function include/fastgeodis_hip.h
  declares; argument validation returns the documented codes;
* the torch glue raises the reference's exception types in the reference's
  order (FastGeodis/common.h:53-120, fastgeodis.cpp:44-146;
  tests/test_fastgeodis.py:46-91, 212-252);
* the Python signatures match the reference wrappers.
"""
import ctypes
import inspect
import os
import re

import pytest
import torch

import fastgeodis_amd as FG

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastgeodis_hip.h")
LIB = os.path.join(ROOT, "fastgeodis_amd", "lib", "libfastgeodis_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fg_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    return ctypes.CDLL(LIB)


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), f"{n} declared in the header but not exported"


def test_abi_version_and_errors(lib):
    lib.fg_abi_version.restype = ctypes.c_int
    assert lib.fg_abi_version() == 4
    lib.fg_last_error.restype = ctypes.c_char_p
    f = lib.fg_geodesic2d_f32
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 4 + [ctypes.c_float] * 3 + \
        [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    rc = f(None, None, None, 1, 1, 8, 8, 1e10, 1.0, 0.0, 2, None, 0, None)
    assert rc == -1
    assert b"null" in lib.fg_last_error()


def test_pending_status_is_read_and_cleared(lib):
    """The per-device status word: a value set from the host is returned once by
    fg_pending_status and then cleared (host logic only, no HIP call)."""
    lib.fg_pending_status.restype = ctypes.c_int
    lib.fg_pending_status.argtypes = [ctypes.c_int]
    lib.fg_inject_status.restype = ctypes.c_int
    lib.fg_inject_status.argtypes = [ctypes.c_int, ctypes.c_int]
    assert lib.fg_pending_status(3) == 0
    assert lib.fg_inject_status(3, -3) == 0
    assert lib.fg_pending_status(2) == 0  # other devices unaffected
    assert lib.fg_pending_status(3) == -3
    assert lib.fg_pending_status(3) == 0
    assert lib.fg_inject_status(-1, -3) == -1
    lib.fg_set_wait_limit.restype = ctypes.c_int
    lib.fg_set_wait_limit.argtypes = [ctypes.c_int]
    assert lib.fg_set_wait_limit(-5) == -1
    assert lib.fg_set_wait_limit(0) == 0


def test_pending_status_through_the_glue():
    FG.FastGeodisCpp.inject_status(5, -3)
    assert FG.FastGeodisCpp.pending_status(5) == -3
    assert FG.FastGeodisCpp.pending_status(5) == 0


def test_workspace_query(lib):
    q = lib.fg_workspace_bytes
    q.restype = ctypes.c_size_t
    q.argtypes = [ctypes.c_int] * 5
    small = q(0, 1, 0, 64, 64)
    big = q(0, 1, 64, 64, 64)
    gsf = q(2, 3, 64, 64, 64)
    assert 0 < small < big < gsf
    # the image transpose and the transposed distances dominate: >= (C+1)*N floats
    assert big >= 2 * 64 ** 3 * 4


def test_default_tuning_is_automatic():
    import threading
    got = {}
    t = threading.Thread(target=lambda: got.update(FG.get_tuning()))  # fresh host thread
    t.start(); t.join()
    assert got == {"engine": 1, "halo": 0, "max_groups": 0, "shape": 0}


def test_auto_halo_workspace_covers_fixed_depths():
    """The automatic halo may pick any depth, so its workspace bounds every fixed one."""
    old = FG.get_tuning()
    try:
        sizes = {}
        for k in range(0, 31):
            FG.set_tuning(1, k, 0, 0)
            sizes[k] = (FG.FastGeodisCpp.workspace_bytes(0, 1, 96, 300, 200),
                        FG.FastGeodisCpp.workspace_bytes(0, 3, 0, 700, 1500))
        assert all(sizes[0][j] >= sizes[k][j] for k in range(1, 31) for j in range(2))
    finally:
        FG.set_tuning(**old)


def test_tuning_roundtrip():
    old = FG.get_tuning()
    try:
        FG.set_tuning(0, 3, 17, 3)
        assert FG.get_tuning() == {"engine": 0, "halo": 3, "max_groups": 17, "shape": 3}
        FG.set_tuning(1, 4, 0, 7)  # the last shape
        assert FG.get_tuning()["shape"] == 7
        with pytest.raises(ValueError):
            FG.set_tuning(1, 4, 0, 8)
        with pytest.raises(ValueError):
            FG.set_tuning(2, 4, 0)
        FG.set_tuning(1, 0, 0)  # automatic halo depth
        assert FG.get_tuning()["halo"] == 0
        with pytest.raises(ValueError):
            FG.set_tuning(1, -1, 0)
        FG.set_tuning(1, 30, 0)  # deepest halo
        assert FG.get_tuning()["halo"] == 30
        with pytest.raises(ValueError):
            FG.set_tuning(1, 31, 0)
    finally:
        FG.set_tuning(**old)


def test_signatures_match_reference_wrappers():
    sig = {n: list(inspect.signature(getattr(FG, n)).parameters) for n in (
        "generalised_geodesic2d", "generalised_geodesic3d", "signed_generalised_geodesic2d",
        "signed_generalised_geodesic3d", "GSF2d", "GSF3d")}
    assert sig["generalised_geodesic2d"] == ["image", "softmask", "v", "lamb", "iter"]
    assert sig["generalised_geodesic3d"] == ["image", "softmask", "spacing", "v", "lamb", "iter"]
    assert sig["signed_generalised_geodesic2d"] == ["image", "softmask", "v", "lamb", "iter"]
    assert sig["signed_generalised_geodesic3d"] == ["image", "softmask", "spacing", "v", "lamb", "iter"]
    assert sig["GSF2d"] == ["image", "softmask", "theta", "v", "lamb", "iter"]
    assert sig["GSF3d"] == ["image", "softmask", "theta", "spacing", "v", "lamb", "iter"]
    assert inspect.signature(FG.generalised_geodesic2d).parameters["iter"].default == 2
    assert inspect.signature(FG.generalised_geodesic3d).parameters["iter"].default == 4


FUNCS = [
    (2, lambda i, m: FG.generalised_geodesic2d(i, m, 1e10, 1.0, 2)),
    (3, lambda i, m: FG.generalised_geodesic3d(i, m, [1.0, 1.0, 1.0], 1e10, 1.0, 2)),
    (2, lambda i, m: FG.signed_generalised_geodesic2d(i, m, 1e10, 1.0, 2)),
    (3, lambda i, m: FG.signed_generalised_geodesic3d(i, m, [1.0, 1.0, 1.0], 1e10, 1.0, 2)),
    (2, lambda i, m: FG.GSF2d(i, m, 0.0, 1e10, 1.0, 2)),
    (3, lambda i, m: FG.GSF3d(i, m, 0.0, [1.0, 1.0, 1.0], 1e10, 1.0, 2)),
]


@pytest.mark.parametrize("nd,fn", FUNCS)
@pytest.mark.parametrize("base", [16, 32])
def test_ill_shape_raises_value_error(nd, fn, base):
    shape = [1, 1] + [base] * nd
    bad_batch = [2] + shape[1:]
    with pytest.raises(ValueError):
        fn(torch.rand(shape), torch.rand(bad_batch))
    with pytest.raises(ValueError):
        fn(torch.rand(bad_batch), torch.rand(bad_batch))
    mism = list(shape)
    mism[-1] = 12
    with pytest.raises(ValueError):
        fn(torch.rand(mism), torch.rand(shape))
    if nd == 2:
        with pytest.raises(ValueError):
            fn(torch.rand(shape + [16]), torch.rand(shape + [16]))


@pytest.mark.parametrize("fn", [
    lambda i, m, s: FG.generalised_geodesic3d(i, m, s, 1e10, 1.0, 2),
    lambda i, m, s: FG.signed_generalised_geodesic3d(i, m, s, 1e10, 1.0, 2),
    lambda i, m, s: FG.GSF3d(i, m, 0.0, s, 1e10, 1.0, 2),
])
def test_ill_spacing_raises_value_error(fn):
    x = torch.zeros(1, 1, 16, 16, 16)
    with pytest.raises(ValueError):
        fn(x, x, [1.0, 1.0])


def test_channel_counts_need_not_match():
    """The reference compares spatial dims only (common.h:53-71): C and Cm may
    differ; the call proceeds past validation (and here stops at the device)."""
    img = torch.rand(1, 3, 8, 8)
    mask = torch.rand(1, 2, 8, 8)
    if torch.cuda.is_available():
        out = FG.generalised_geodesic2d(img, mask, 1e10, 1.0, 2)
        assert out.shape == mask.shape
    else:
        with pytest.raises(RuntimeError):
            FG.generalised_geodesic2d(img, mask, 1e10, 1.0, 2)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_cpu_tensors_without_gpu_raise_runtime_error():
    x = torch.zeros(1, 1, 8, 8)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        FG.generalised_geodesic2d(x, x, 1e10, 1.0, 2)


def test_native_libraries_are_in_tree():
    for p in FG.native_library_paths():
        assert os.path.isfile(p)
        assert os.path.commonpath([p, ROOT]) == ROOT
