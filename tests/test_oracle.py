"""Pins the CPU oracle (oracle/fg_oracle.c) before it is trusted as the checker.

* golden fixtures written by the reference's own CPU extension (bit-exact);
* the reference build itself on fresh seeded inputs, when it is present;
* the float32 Dijkstra fixpoint at convergence (bit-exact);
* closed forms: iter 0, zero/one masks, lamb=1 on a constant image,
  Euclidean chamfer distances with lamb=0.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from oracle.dijkstra import geodesic_fixpoint

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load_golden():
    with open(os.path.join(GOLDEN, "cases.json")) as f:
        cases = json.load(f)
    data = np.load(os.path.join(GOLDEN, "fixtures.npz"))
    return cases, data


def run_oracle_case(case, image, mask):
    v, lamb, it = case["v"], case["lamb"], case["iter"]
    nd, kind = case["nd"], case["kind"]
    sp = case["spacing"]
    if kind == "unsigned":
        return oracle.py_geodesic2d(image, mask, v, lamb, it) if nd == 2 else \
            oracle.py_geodesic3d(image, mask, sp, v, lamb, it)
    if kind == "signed":
        return oracle.py_signed_geodesic2d(image, mask, v, lamb, it) if nd == 2 else \
            oracle.py_signed_geodesic3d(image, mask, sp, v, lamb, it)
    return oracle.py_gsf2d(image, mask, case["theta"], v, lamb, it) if nd == 2 else \
        oracle.py_gsf3d(image, mask, case["theta"], sp, v, lamb, it)


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


CASES, DATA = load_golden()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] + "-" + c["kind"] for c in CASES])
def test_oracle_matches_golden(case):
    n = case["name"]
    out = run_oracle_case(case, DATA[n + "_image"], DATA[n + "_mask"])
    assert bits_equal(out, DATA[n + "_out"]), f"{n}: max |diff| {np.nanmax(np.abs(out - DATA[n + '_out']))}"


def test_golden_covers_the_path():
    kinds = {(c["kind"], c["nd"]) for c in CASES}
    for k in ("unsigned", "signed", "gsf"):
        for nd in (2, 3):
            assert (k, nd) in kinds
    assert any(c["C"] > 1 for c in CASES) and any(c["Cm"] > 1 for c in CASES)
    assert any(c["iter"] == 0 for c in CASES)
    assert any(c["mask"] == "seeds_nan" for c in CASES)
    assert any(c["nd"] == 3 and c["spacing"] and len(set(c["spacing"])) == 3 for c in CASES)


@pytest.mark.skipif(oracle.reference_module() is None, reason="reference CPU build not present")
@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_reference_build(seed):
    import torch

    ref = oracle.reference_module()
    rng = np.random.default_rng(100 + seed)
    for nd in (2, 3):
        C = int(rng.integers(1, 4))
        shape = [int(x) for x in rng.integers(2, 40 if nd == 2 else 14, size=nd)]
        img = rng.random((1, C, *shape), dtype=np.float32)
        mask = (rng.random((1, 1, *shape)) > 0.03).astype(np.float32)
        lamb = float(rng.choice([0.0, 0.2, 0.5, 1.0]))
        it = int(rng.integers(1, 4))
        sp = [float(x) for x in rng.uniform(0.4, 3.0, size=3)]
        ti, tm = torch.from_numpy(img), torch.from_numpy(mask)
        if nd == 2:
            r = ref.signed_generalised_geodesic2d(ti, tm, 1e10, lamb, 1 - lamb, it).numpy()
            o = oracle.py_signed_geodesic2d(img, mask, 1e10, lamb, it)
        else:
            r = ref.signed_generalised_geodesic3d(ti, tm, sp, 1e10, lamb, 1 - lamb, it).numpy()
            o = oracle.py_signed_geodesic3d(img, mask, sp, 1e10, lamb, it)
        assert bits_equal(o, r)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_converges_to_dijkstra(seed):
    rng = np.random.default_rng(seed)
    nd = 2 if seed % 2 == 0 else 3
    C = int(rng.integers(1, 4))
    shape = [int(x) for x in rng.integers(3, 12 if nd == 2 else 7, size=nd)]
    img = rng.random((1, C, *shape), dtype=np.float32)
    mask = np.ones((1, 1, *shape), np.float32)
    for _ in range(2):
        mask[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
    lamb = float(rng.choice([0.0, 0.5, 0.9, 1.0]))
    lg, le = oracle.wrapper_weights(lamb)
    sp = [float(x) for x in rng.uniform(0.5, 3.0, size=3)] if nd == 3 else None
    prev = None
    for it in range(1, 80):
        out = oracle.geodesic2d(img, mask, 1e10, lg, le, it) if nd == 2 else \
            oracle.geodesic3d(img, mask, sp, 1e10, lg, le, it)
        if prev is not None and bits_equal(prev, out):
            break
        prev = out
    ref = geodesic_fixpoint(img, mask, 1e10, lg, le, sp)
    assert bits_equal(out[0, 0], ref)


def test_dijkstra_detects_a_wrong_table():
    """A planted mistake in the oracle's 3D tables (W pass summed in the D
    pass order) must change the converged result for anisotropic spacing."""
    rng = np.random.default_rng(7)
    shape = (5, 6, 7)
    img = np.zeros((1, 1, *shape), np.float32)
    mask = np.ones((1, 1, *shape), np.float32)
    mask[0, 0, 0, 0, 0] = 0
    sp = [1.7, 0.3, 2.9]
    lg, le = oracle.wrapper_weights(0.0)
    out = None
    for it in range(1, 20):
        o = oracle.geodesic3d(img, mask, sp, 1e10, lg, le, it)
        if out is not None and bits_equal(o, out):
            break
        out = o
    good = geodesic_fixpoint(img, mask, 1e10, lg, le, sp)
    assert bits_equal(out[0, 0], good)
    # wrong spacing-to-axis mapping must be visible
    bad = geodesic_fixpoint(img, mask, 1e10, lg, le, [sp[2], sp[1], sp[0]])
    assert not bits_equal(out[0, 0], bad)


# ---- closed forms -----------------------------------------------------------

def test_iter_zero_returns_v_times_mask():
    rng = np.random.default_rng(1)
    img = rng.random((1, 2, 9, 11), dtype=np.float32)
    mask = rng.random((1, 2, 9, 11), dtype=np.float32)
    out = oracle.py_geodesic2d(img, mask, 37.5, 0.5, 0)
    assert bits_equal(out, np.float32(37.5) * mask)


@pytest.mark.parametrize("nd", [2, 3])
def test_zero_and_one_masks(nd):
    shape = (1, 1) + (8,) * nd
    z = np.zeros(shape, np.float32)
    o = np.ones(shape, np.float32)
    f = (lambda i, m: oracle.py_geodesic2d(i, m, 1e10, 1.0, 2)) if nd == 2 else \
        (lambda i, m: oracle.py_geodesic3d(i, m, [1, 1, 1], 1e10, 1.0, 2))
    assert np.all(f(z, z) == 0)
    assert np.all(f(z, o) == np.float32(1e10))
    s = (lambda i, m: oracle.py_signed_geodesic2d(i, m, 1e10, 1.0, 2)) if nd == 2 else \
        (lambda i, m: oracle.py_signed_geodesic3d(i, m, [1, 1, 1], 1e10, 1.0, 2))
    assert np.all(s(z, z) == np.float32(-1e10))
    assert np.all(s(z, o) == np.float32(1e10))


@pytest.mark.parametrize("nd", [2, 3])
def test_lamb_one_constant_image_is_zero_after_one_iteration(nd):
    rng = np.random.default_rng(2)
    shape = (1, 1) + ((13, 17) if nd == 2 else (5, 7, 6))
    img = np.full(shape, 0.3, np.float32)
    mask = np.ones(shape, np.float32)
    mask[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape[2:])] = 0
    out = oracle.py_geodesic2d(img, mask, 1e10, 1.0, 1) if nd == 2 else \
        oracle.py_geodesic3d(img, mask, [2.0, 1.0, 0.5], 1e10, 1.0, 1)
    assert np.all(out == 0)


def test_euclidean_chamfer_2d():
    """lamb=0, uniform image, corner seed: one iteration gives the 8-neighbour
    chamfer distance (axis steps 1, diagonal steps sqrt 2)."""
    H, W = 19, 23
    img = np.ones((1, 1, H, W), np.float32)
    mask = np.ones((1, 1, H, W), np.float32)
    mask[0, 0, 0, 0] = 0
    out = oracle.py_geodesic2d(img, mask, 1e10, 0.0, 1)[0, 0]
    h, w = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    lo, hi = np.minimum(h, w), np.maximum(h, w)
    exact = (hi - lo) + math.sqrt(2) * lo
    np.testing.assert_allclose(out, exact, rtol=1e-5, atol=1e-5)


def test_reference_euclidean_band():
    """Reference tests/test_fastgeodis.py:162-210: max distance within +-10 of
    sqrt(n)*b for a corner seed with lamb=0."""
    for nd, b in ((2, 32), (3, 16)):
        shape = (1, 1) + (b,) * nd
        img = np.ones(shape, np.float32)
        mask = np.ones(shape, np.float32)
        mask[(0, 0) + (0,) * nd] = 0
        out = oracle.py_geodesic2d(img, mask, 1e10, 0.0, 2) if nd == 2 else \
            oracle.py_geodesic3d(img, mask, [1, 1, 1], 1e10, 0.0, 2)
        exp = math.sqrt(nd * b * b)
        tol = 10 if nd == 2 else 100
        assert exp - tol < out.max() < exp + tol


@pytest.mark.parametrize("seed", range(4))
def test_layout_faithful_forms_equal_the_oracle(seed):
    """The CPU-baseline forms (per-pass contiguous copies like the reference's
    transposes) return the oracle's bits."""
    rng = np.random.default_rng(300 + seed)
    C = int(rng.integers(1, 4))
    lamb = float(rng.choice([0.0, 0.5, 1.0]))
    lg, le = oracle.wrapper_weights(lamb)
    it = int(rng.integers(1, 4))
    shape2 = [int(x) for x in rng.integers(1, 70, size=2)]
    img = rng.random((1, C, *shape2), dtype=np.float32)
    mask = (rng.random((1, 1, *shape2)) > 0.05).astype(np.float32)
    assert bits_equal(oracle.geodesic2d_layout(img, mask, 1e10, lg, le, it),
                      oracle.geodesic2d(img, mask, 1e10, lg, le, it))
    shape3 = [int(x) for x in rng.integers(1, 30, size=3)]
    sp = [float(x) for x in rng.uniform(0.4, 3.0, size=3)]
    img = rng.random((1, C, *shape3), dtype=np.float32)
    mask = (rng.random((1, 1, *shape3)) > 0.05).astype(np.float32)
    assert bits_equal(oracle.geodesic3d_layout(img, mask, sp, 1e10, lg, le, it),
                      oracle.geodesic3d(img, mask, sp, 1e10, lg, le, it))
