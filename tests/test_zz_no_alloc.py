"""The no-allocation entry point (fg_geodesic3d_f32_noalloc) against the
workspace entry point (fg_geodesic3d_f32) through the C ABI: bit-identical
results on two calls in a row, the second one smaller (the shared scratch keeps
the first call's contents), and a clear error for a volume beyond the scratch.
"""
import ctypes

import numpy as np
import pytest
import torch

import fastgeodis_amd as FG

pytestmark = pytest.mark.gpu


def lib_funcs():
    lib = ctypes.CDLL(FG.native_library_paths()[0])
    ws_f = lib.fg_geodesic3d_f32
    ws_f.restype = ctypes.c_int
    ws_f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_float)] + \
        [ctypes.c_float] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    na_f = lib.fg_geodesic3d_f32_noalloc
    na_f.restype = ctypes.c_int
    na_f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_float)] + \
        [ctypes.c_float] * 3 + [ctypes.c_int, ctypes.c_void_p]
    q = lib.fg_workspace_bytes
    q.restype = ctypes.c_size_t
    q.argtypes = [ctypes.c_int] * 5
    cap = lib.fg_noalloc_scratch_bytes
    cap.restype = ctypes.c_size_t
    st = lib.fg_status
    st.restype = ctypes.c_int
    st.argtypes = [ctypes.c_void_p]
    err = lib.fg_last_error
    err.restype = ctypes.c_char_p
    return ws_f, na_f, q, cap, st, err


def case(seed, shape, C):
    rng = np.random.default_rng(seed)
    img = torch.from_numpy(rng.random((1, C, *shape), dtype=np.float32)).cuda()
    mask = torch.ones((1, 1, *shape), dtype=torch.float32, device="cuda")
    for _ in range(5):
        mask[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
    return img, mask


def test_noalloc_equals_workspace_entry_point(device):
    ws_f, na_f, q, cap, st, err = lib_funcs()
    stream = torch.cuda.current_stream().cuda_stream
    for k, (shape, C, sp, lamb, it) in enumerate([((70, 200, 150), 2, (2.0, 1.0, 0.5), 0.4, 3),
                                                   ((33, 41, 90), 1, (1.0, 1.0, 1.0), 1.0, 2)]):
        print(f"[no_alloc] call {k}: shape {shape} C={C}", flush=True)
        img, mask = case(10 + k, shape, C)
        spa = (ctypes.c_float * 3)(*sp)
        lg, le = np.float32(lamb), np.float32(1.0 - lamb)
        want = torch.empty_like(mask)
        wsb = q(0, C, *shape)
        ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
        assert ws_f(img.data_ptr(), mask.data_ptr(), want.data_ptr(), C, 1, *shape, spa, 1e10, float(lg),
                    float(le), it, ws.data_ptr(), wsb, stream) == 0
        got = torch.full_like(mask, float("nan"))
        rc = na_f(img.data_ptr(), mask.data_ptr(), got.data_ptr(), C, 1, *shape, spa, 1e10, float(lg), float(le),
                  it, stream)
        assert rc == 0, err()
        assert st(stream) == 0, err()
        assert torch.equal(got.view(torch.int32), want.view(torch.int32)), f"call {k} differs"


def test_noalloc_rejects_volumes_beyond_its_scratch(device):
    ws_f, na_f, q, cap, st, err = lib_funcs()
    assert cap() >= 200 << 20
    # a volume whose workspace exceeds the scratch: refused before anything is
    # launched (the pointers are never dereferenced)
    x = torch.zeros(16, device="cuda")
    spa = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    rc = na_f(x.data_ptr(), x.data_ptr(), x.data_ptr(), 1, 1, 512, 512, 512, spa, 1e10, 1.0, 0.0, 2,
              torch.cuda.current_stream().cuda_stream)
    assert rc == -4
    assert b"no-allocation" in err()
