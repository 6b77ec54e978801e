"""GPU parity: the HIP path against the oracle and the reference's golden
vectors, bit for bit (FP32, no FMA, the reference's neighbour order).

Every case goes through the reference-compatible API (FastGeodisCpp glue ->
C ABI -> HIP kernels) on cuda:0.  FASTGEODIS_SYNC_CHECK=1 (set in conftest)
makes each call synchronise and check the device status word, so an expired
halo wait surfaces as RuntimeError.
"""
import json
import os
import sys

import numpy as np
import pytest
import torch

import fastgeodis_amd as FG
import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "cases.json")) as _f:
    CASES = json.load(_f)
DATA = np.load(os.path.join(HERE, "golden", "fixtures.npz"))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def diff_report(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    idx = np.argwhere(bad)
    return f"{bad.sum()} mismatches; first {idx[:3].tolist()} got {a[bad][:3]} want {b[bad][:3]}"


# (engine, halo K, shape); K is rounded up to a multiple of the shape's slab
# ring (4x8: 2, 8x4: 3, 16x2: 3, 8x4 deep: 5, 16x2 deep: 5, 8x4 deep with
# exchange waves: 5, 8x2 deep: 5; shape 0 = 16x2 deep; one-row planes: 6, or
# 30 for K > 24), K = 0 picks it per plane
ENGINES = [(1, 0, 0), (1, 4, 0), (1, 1, 1), (1, 3, 3), (1, 7, 2), (1, 2, 1), (1, 5, 3), (1, 5, 4),
           (1, 10, 4), (1, 0, 4), (1, 10, 5), (0, 4, 0), (1, 0, 6), (1, 10, 6), (1, 24, 0), (1, 30, 0),
           (1, 0, 7), (1, 5, 7)]


@pytest.fixture(params=ENGINES, ids=["tiled-auto", "tiled-K4-auto", "tiled-K1-4x8", "tiled-K3-16x2",
                                     "tiled-K7-8x4", "tiled-K2-4x8", "tiled-K5-16x2", "tiled-K5-8x4deep",
                                     "tiled-K10-8x4deep", "tiled-auto-8x4deep", "tiled-K10-16x2deep", "per-slice",
                                     "tiled-auto-8x4xchg", "tiled-K10-8x4xchg", "tiled-K24", "tiled-K30",
                                     "tiled-auto-8x2", "tiled-K5-8x2"])
def engine(request, device):
    e, k, shape = request.param
    old = FG.get_tuning()
    FG.set_tuning(e, k, 0, shape)
    yield request.param
    FG.set_tuning(**old)


def run_api(case, image, mask):
    dev = "cuda"
    ti = torch.from_numpy(image).to(dev)
    tm = torch.from_numpy(mask).to(dev)
    v, lamb, it, nd, kind = case["v"], case["lamb"], case["iter"], case["nd"], case["kind"]
    sp = case.get("spacing")
    if kind == "unsigned":
        r = FG.generalised_geodesic2d(ti, tm, v, lamb, it) if nd == 2 else \
            FG.generalised_geodesic3d(ti, tm, sp, v, lamb, it)
    elif kind == "signed":
        r = FG.signed_generalised_geodesic2d(ti, tm, v, lamb, it) if nd == 2 else \
            FG.signed_generalised_geodesic3d(ti, tm, sp, v, lamb, it)
    else:
        r = FG.GSF2d(ti, tm, case["theta"], v, lamb, it) if nd == 2 else \
            FG.GSF3d(ti, tm, case["theta"], sp, v, lamb, it)
    torch.cuda.synchronize()
    return r.cpu().numpy()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] + "-" + c["kind"] for c in CASES])
def test_golden(case, engine):
    n = case["name"]
    say(f"golden {n} {case['kind']} nd={case['nd']} C={case['C']} shape={case['shape']} engine={engine}")
    out = run_api(case, DATA[n + "_image"], DATA[n + "_mask"])
    want = DATA[n + "_out"]
    assert bits_equal(out, want), diff_report(out, want)


def blobs(rng, shape, nblobs=6):
    grids = np.meshgrid(*[np.arange(s, dtype=np.float32) for s in shape], indexing="ij")
    f = np.zeros(shape, np.float32)
    for _ in range(nblobs):
        c = [rng.uniform(0, s) for s in shape]
        sig = rng.uniform(2.0, max(3.0, max(shape) / 3))
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        f += np.float32(rng.uniform(0, 1)) * np.exp(-r2 / (2 * sig * sig)).astype(np.float32)
    return f + rng.uniform(0, 0.05, size=shape).astype(np.float32)


# shapes chosen so that every pass of the tiled engine uses several tiles
# along both in-plane axes (E > 62 lanes, B > 32 rows) and partial last tiles
RANDOM_3D = [
    ((37, 70, 90), [2.5, 1.0, 1.0], 1, 1.0, 4),
    ((67, 41, 129), [0.7, 1.9, 1.3], 1, 0.5, 2),
    ((33, 65, 66), [1.0, 1.0, 1.0], 3, 0.0, 2),
    ((20, 100, 64), [1.2, 1.0, 3.0], 2, 0.8, 3),
    ((9, 35, 140), [1.0, 2.0, 1.0], 4, 1.0, 2),
    ((70, 8, 70), [1.0, 1.0, 1.0], 1, 0.3, 2),
    ((5, 130, 33), [3.0, 0.5, 1.0], 5, 0.6, 2),      # C=5: per-slice fallback
    ((40, 40, 40), [1.0, 1.0, 1.0], 1, 1.25, 2),     # lamb > 1: exact-order path
    ((31, 47, 71), [1.0, 1.0, 1.0], 1, -0.5, 2),     # negative lamb: exact-order path
    # degenerate sides: a pass with one slice does nothing, planes of one or
    # two rows/lanes, a single voxel
    ((1, 50, 70), [1.0, 1.0, 1.0], 1, 1.0, 2),
    ((30, 1, 90), [2.0, 1.0, 1.0], 2, 0.5, 2),
    ((25, 40, 1), [1.0, 1.0, 1.0], 1, 0.0, 3),
    ((2, 2, 2), [1.0, 1.0, 1.0], 1, 1.0, 2),
    ((1, 1, 1), [1.0, 1.0, 1.0], 1, 0.5, 2),
    ((3, 200, 2), [1.0, 0.5, 2.0], 1, 0.7, 2),
    ((64, 2, 300), [1.5, 1.0, 1.0], 3, 1.0, 2),
    # W-backward sweeps writing the native layout in runs of 16 slices: a
    # partial first run (W - 1 = 16, 32), W below one run, several B tiles
    ((90, 33, 17), [1.0, 1.3, 0.8], 1, 1.0, 2),
    ((70, 66, 33), [2.0, 1.0, 1.0], 1, 0.5, 3),
    ((130, 40, 9), [1.0, 1.0, 1.0], 1, 0.0, 2),
]


@pytest.mark.parametrize("shape,spacing,C,lamb,it", RANDOM_3D)
def test_random_3d_vs_oracle(shape, spacing, C, lamb, it, engine):
    rng = np.random.default_rng(hash((shape, C)) % 2**31)
    img = np.stack([blobs(rng, shape) for _ in range(C)])[None]
    mask = np.ones((1, 1, *shape), np.float32)
    for _ in range(5):
        mask[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
    say(f"3d {shape} C={C} lamb={lamb} it={it} engine={engine}")
    want = oracle.py_geodesic3d(img, mask, spacing, 1e10, lamb, it)
    out = FG.generalised_geodesic3d(torch.from_numpy(img).cuda(), torch.from_numpy(mask).cuda(),
                                    spacing, 1e10, lamb, it).cpu().numpy()
    assert bits_equal(out, want), diff_report(out, want)


RANDOM_2D = [((300, 257), 1, 1.0, 2), ((61, 62), 3, 0.5, 2), ((1, 700), 1, 0.7, 3),
             ((513, 1), 2, 0.2, 2), ((128, 1000), 3, 0.0, 1), ((97, 130), 4, 1.0, 4),
             # row and column sweeps of exactly one 30-slice block, and of two
             ((31, 200), 1, 1.0, 2), ((61, 31), 2, 0.5, 3)]


@pytest.mark.parametrize("shape,C,lamb,it", RANDOM_2D)
def test_random_2d_vs_oracle(shape, C, lamb, it, engine):
    rng = np.random.default_rng(sum(shape) + C)
    img = np.stack([blobs(rng, shape) for _ in range(C)])[None]
    mask = np.ones((1, 1, *shape), np.float32)
    for _ in range(4):
        mask[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
    say(f"2d {shape} C={C} lamb={lamb} it={it} engine={engine}")
    want = oracle.py_geodesic2d(img, mask, 1e10, lamb, it)
    out = FG.generalised_geodesic2d(torch.from_numpy(img).cuda(), torch.from_numpy(mask).cuda(),
                                    1e10, lamb, it).cpu().numpy()
    assert bits_equal(out, want), diff_report(out, want)


@pytest.mark.parametrize("nd", [2, 3])
def test_signed_and_gsf_vs_oracle(nd, device):
    rng = np.random.default_rng(11 + nd)
    shape = (150, 97) if nd == 2 else (23, 45, 80)
    sp = [1.5, 1.0, 0.8]
    img = np.stack([blobs(rng, shape) for _ in range(3)])[None]
    f = blobs(rng, shape)
    f = (f - f.min()) / (f.max() - f.min())
    mask = (f > 0.5).astype(np.float32)[None, None]
    ti, tm = torch.from_numpy(img).cuda(), torch.from_numpy(mask).cuda()
    say(f"signed/gsf nd={nd}")
    if nd == 2:
        want = oracle.py_signed_geodesic2d(img, mask, 1e10, 0.5, 2)
        out = FG.signed_generalised_geodesic2d(ti, tm, 1e10, 0.5, 2).cpu().numpy()
        assert bits_equal(out, want), diff_report(out, want)
        want = oracle.py_gsf2d(img, mask, 0.1, 1e10, 1.0, 4)
        out = FG.GSF2d(ti, tm, 0.1, 1e10, 1.0, 4).cpu().numpy()
    else:
        want = oracle.py_signed_geodesic3d(img, mask, sp, 1e10, 0.5, 2)
        out = FG.signed_generalised_geodesic3d(ti, tm, sp, 1e10, 0.5, 2).cpu().numpy()
        assert bits_equal(out, want), diff_report(out, want)
        want = oracle.py_gsf3d(img, mask, 0.1, sp, 1e10, 1.0, 2)
        out = FG.GSF3d(ti, tm, 0.1, sp, 1e10, 1.0, 2).cpu().numpy()
    assert bits_equal(out, want), diff_report(out, want)


# ---- the reference's own invariants (tests/test_fastgeodis.py) on the GPU ----

@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("base", [16, 32, 64])
def test_reference_invariants(nd, base, device):
    shape = [1, 1] + [base] * nd
    f = (lambda i, m, l: FG.generalised_geodesic2d(i, m, 1e10, l, 2)) if nd == 2 else \
        (lambda i, m, l: FG.generalised_geodesic3d(i, m, [1.0, 1.0, 1.0], 1e10, l, 2))
    s = (lambda i, m: FG.signed_generalised_geodesic2d(i, m, 1e10, 1.0, 2)) if nd == 2 else \
        (lambda i, m: FG.signed_generalised_geodesic3d(i, m, [1.0, 1.0, 1.0], 1e10, 1.0, 2))
    z = torch.zeros(shape, device="cuda")
    o = torch.ones(shape, device="cuda")
    say(f"invariants nd={nd} base={base}")
    assert torch.all(f(z, z, 1.0) == 0)
    assert torch.all(f(z, o, 1.0) == np.float32(1e10))
    assert torch.all(s(z, z) == np.float32(-1e10))
    assert torch.all(s(z, o) == np.float32(1e10))
    m = o.clone()
    m[(0, 0) + (0,) * nd] = 0
    d = f(o, m, 0.0)
    exp = (nd * base * base) ** 0.5
    tol = 10 if nd == 2 else 100
    assert exp - tol < float(d.max()) < exp + tol
    # GSF runs on the same shapes (reference tests/test_fastgeodis.py:416-436)
    g = FG.GSF2d(torch.rand(shape, device="cuda"), torch.rand(shape, device="cuda"), 0.0, 1e10, 1.0, 2) \
        if nd == 2 else FG.GSF3d(torch.rand(shape, device="cuda"), torch.rand(shape, device="cuda"), 0.0,
                                 [1.0, 1.0, 1.0], 1e10, 1.0, 2)
    assert g.shape == tuple(shape)


def test_cpu_tensors_are_staged_through_the_gpu(device):
    rng = np.random.default_rng(5)
    img = rng.random((1, 1, 40, 50), dtype=np.float32)
    mask = (rng.random((1, 1, 40, 50)) > 0.02).astype(np.float32)
    out = FG.generalised_geodesic2d(torch.from_numpy(img), torch.from_numpy(mask), 1e10, 1.0, 2)
    assert out.device.type == "cpu"
    assert bits_equal(out.numpy(), oracle.py_geodesic2d(img, mask, 1e10, 1.0, 2))


def test_device_mismatch_raises_value_error(device):
    x = torch.rand(1, 1, 16, 16)
    with pytest.raises(ValueError):
        FG.generalised_geodesic2d(x.cuda(), x, 1e10, 1.0, 2)
    with pytest.raises(ValueError):
        FG.generalised_geodesic2d(x, x.cuda(), 1e10, 1.0, 2)


def test_non_float_dtype_raises_runtime_error(device):
    x = torch.rand(1, 1, 16, 16, device="cuda")
    with pytest.raises(RuntimeError):
        FG.generalised_geodesic2d(x.double(), x, 1e10, 1.0, 2)


def test_repeat_runs_are_bit_identical(device):
    rng = np.random.default_rng(9)
    shape = (64, 96, 130)
    img = torch.from_numpy(blobs(rng, shape)[None, None]).cuda()
    mask = torch.ones_like(img)
    mask[0, 0, 10, 20, 30] = 0
    say("repeat 64x96x130")
    first = FG.generalised_geodesic3d(img, mask, [1.0, 1.0, 1.0], 1e10, 1.0, 2)
    for _ in range(5):
        again = FG.generalised_geodesic3d(img, mask, [1.0, 1.0, 1.0], 1e10, 1.0, 2)
        assert torch.equal(first.view(torch.int32), again.view(torch.int32))


def test_c_abi_direct(device):
    """Calls libfastgeodis_hip.so through ctypes with torch-owned buffers."""
    import ctypes

    lib = ctypes.CDLL(FG.native_library_paths()[0])
    f = lib.fg_geodesic3d_f32
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_float)] + \
        [ctypes.c_float] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    q = lib.fg_workspace_bytes
    q.restype = ctypes.c_size_t
    q.argtypes = [ctypes.c_int] * 5
    st = lib.fg_status
    st.restype = ctypes.c_int
    st.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(3)
    shape = (21, 66, 70)
    img = rng.random((1, 2, *shape), dtype=np.float32)
    mask = np.ones((1, 1, *shape), np.float32)
    mask[0, 0, 3, 4, 5] = 0
    ti, tm = torch.from_numpy(img).cuda(), torch.from_numpy(mask).cuda()
    out = torch.empty_like(tm)
    wsb = q(0, 2, *shape)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    sp = (ctypes.c_float * 3)(2.0, 1.0, 0.5)
    stream = torch.cuda.current_stream().cuda_stream
    lg, le = oracle.wrapper_weights(0.4)
    rc = f(ti.data_ptr(), tm.data_ptr(), out.data_ptr(), 2, 1, *shape, sp, 1e10, float(lg), float(le), 3,
           ws.data_ptr(), wsb, stream)
    assert rc == 0
    assert st(stream) == 0
    want = oracle.geodesic3d(img, mask, [2.0, 1.0, 0.5], 1e10, lg, le, 3)
    assert bits_equal(out.cpu().numpy(), want)
